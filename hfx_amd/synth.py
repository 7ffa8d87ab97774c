"""Seeded synthetic workloads (proteins and MS/MS spectra) for the search path.

Proteins: lengths uniform in [min_len, max_len], residues drawn i.i.d. from the
UniProtKB/Swiss-Prot background frequencies.  Spectra: a random database
peptide at a chosen precursor charge; singly charged b and y ions (and doubly
charged ones for 3+ precursors when requested) with intensity U(0.2, 1), plus
three times as many uniform noise peaks (m/z U(50, 2000), intensity U(0, 0.3)),
all m/z <= 2000, sorted by m/z.  Precursor m/z = (M + z*H)/z with M the
peptide's neutral mass (host constants of Mass.hs:19-27).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

AA = "ARNDCQEGHILKMFPSTWYV"
# UniProtKB/Swiss-Prot background frequencies (%), same order as AA
FREQ = np.array(
    [8.25, 5.53, 4.06, 5.45, 1.37, 3.93, 6.75, 7.07, 2.27, 5.96,
     9.66, 5.84, 2.42, 3.86, 4.70, 6.56, 5.34, 1.08, 2.92, 6.87]
)
H2O = 18.010565
PROTON = 1.007276
# monoisotopic residue masses (Mass.hs:147-215), letters A..Z
MONO = {
    "A": 71.037114, "B": 114.53494, "C": 103.009185, "D": 115.026943, "E": 129.042593,
    "F": 147.068414, "G": 57.021464, "H": 137.058912, "I": 113.084064, "J": 113.16472,
    "K": 128.094963, "L": 113.084064, "M": 131.040485, "N": 114.042927, "O": 114.07931,
    "P": 97.052764, "Q": 128.058578, "R": 156.101111, "S": 87.032028, "T": 101.047679,
    "U": 150.04344, "V": 99.068414, "W": 186.079313, "X": 113.08406, "Y": 163.06332,
    "Z": 128.55059,
}


@dataclass
class Proteins:
    seq: np.ndarray  # uint8 letters, all proteins concatenated
    seg: np.ndarray  # uint32 [n+1] offsets

    @property
    def count(self) -> int:
        return len(self.seg) - 1


def make_proteins(n: int, seed: int, min_len: int = 100, max_len: int = 800) -> Proteins:
    rng = np.random.default_rng(seed)
    lens = rng.integers(min_len, max_len + 1, size=n)
    seg = np.zeros(n + 1, dtype=np.uint32)
    np.cumsum(lens, out=seg[1:])
    p = FREQ / FREQ.sum()
    letters = np.frombuffer(AA.encode(), dtype=np.uint8)
    seq = letters[rng.choice(len(AA), size=int(seg[-1]), p=p)]
    return Proteins(seq=np.ascontiguousarray(seq), seg=seg)


@dataclass
class Spectra:
    pcr: np.ndarray          # float32 precursor m/z
    charge: np.ndarray       # float32 precursor charge
    mz: np.ndarray           # float32 all peaks concatenated (sorted per spectrum)
    inten: np.ndarray        # float32
    off: np.ndarray          # uint32 [n+1]
    source: np.ndarray       # int64 database peptide each spectrum was made from
    info: dict = field(default_factory=dict)

    @property
    def count(self) -> int:
        return len(self.pcr)


def peptide_residues(prot: Proteins, tc: int, tn: int) -> bytes:
    """Residues of a digested peptide: [tc, tn] for cleaved pieces, [tc, tn)
    for a protein's final piece (tn == protein end)."""
    p = int(np.searchsorted(prot.seg, tc, side="right")) - 1
    end = int(prot.seg[p + 1])
    last = tn if tn == end else tn + 1
    return prot.seq[tc:last].tobytes()


def make_spectra(prot: Proteins, r: np.ndarray, tc: np.ndarray, tn: np.ndarray, n: int,
                 seed: int, charges=(2,), charge_p=None, static_mods: dict | None = None,
                 double_for_3plus: bool = False, mz_max: float = 2000.0,
                 var_mods: dict | None = None, max_sites: int = 3,
                 mod_p: float = 0.5) -> Spectra:
    """var_mods: with probability mod_p a spectrum's peptide carries 1..max_sites
    variable modifications (residue -> mass delta) at random modifiable
    residues (drawn from a second generator, so the rest of the workload is
    the same with or without them)."""
    rng = np.random.default_rng(seed)
    mrng = np.random.default_rng(seed + 7919)
    masses = dict(MONO)
    for k, v in (static_mods or {}).items():
        masses[k] = masses[k] + v
    src = rng.integers(0, len(r), size=n)
    ch = rng.choice(np.asarray(charges), size=n, p=charge_p)
    pcr = np.empty(n, dtype=np.float32)
    mz_all, in_all = [], []
    off = np.zeros(n + 1, dtype=np.uint32)
    nmod = np.zeros(n, np.int32)
    for i in range(n):
        res = peptide_residues(prot, int(tc[src[i]]), int(tn[src[i]]))
        m = np.array([masses[chr(c)] for c in res])
        if var_mods and mrng.random() < mod_p:
            sites = [k for k, c in enumerate(res) if chr(c) in var_mods]
            if sites:
                k = int(mrng.integers(1, min(max_sites, len(sites)) + 1))
                for q in mrng.choice(sites, size=k, replace=False):
                    m[q] += var_mods[chr(res[q])]
                nmod[i] = k
        b = np.cumsum(m)[:-1]
        tot = float(m.sum())
        y = tot - b
        z = int(ch[i])
        neutral = tot + H2O
        pcr[i] = np.float32((neutral + z * PROTON) / z)
        ions = [b + PROTON, y + H2O + PROTON]
        if double_for_3plus and z >= 3:
            ions += [(b + 2 * PROTON) / 2, (y + H2O + 2 * PROTON) / 2]
        sig = np.concatenate(ions)
        sig = sig[sig <= mz_max]
        nsig = len(sig)
        noise = rng.uniform(50.0, mz_max, size=3 * max(nsig, 1))
        mz = np.concatenate([sig, noise])
        it = np.concatenate([rng.uniform(0.2, 1.0, size=nsig), rng.uniform(0.0, 0.3, size=len(noise))])
        o = np.argsort(mz, kind="stable")
        mz_all.append(mz[o].astype(np.float32))
        in_all.append(it[o].astype(np.float32))
        off[i + 1] = off[i] + len(mz)
    return Spectra(
        pcr=pcr,
        charge=ch.astype(np.float32),
        mz=np.concatenate(mz_all) if mz_all else np.zeros(0, np.float32),
        inten=np.concatenate(in_all) if in_all else np.zeros(0, np.float32),
        off=off,
        source=src.astype(np.int64),
        info={"mod_sites": nmod} if var_mods else {},
    )


# Workloads named by what they are (sizes as in the benchmark definitions).
WORKLOADS = {
    "tryptic-2k-100": dict(
        proteins=2000, spectra=100, seed=1, missed=1, p_rule=1, min_len=7, max_len=30,
        charges=(2,), charge_p=None, static={}, double3=False, topn=5,
    ),
    "tryptic-20k-10k": dict(
        proteins=20000, spectra=10000, seed=2, missed=2, p_rule=0, min_len=0, max_len=0,
        charges=(2, 3), charge_p=(0.6, 0.4), static={"C": 57.02146}, double3=True, topn=10,
    ),
    # the dense-block workload: the same database, 20,000 spectra at 2+/3+,
    # no variable modifications, top-5
    "tryptic-20k-20k": dict(
        proteins=20000, spectra=20000, seed=4, missed=2, p_rule=0, min_len=0, max_len=0,
        charges=(2, 3), charge_p=(0.5, 0.5), static={"C": 57.02146}, double3=True, topn=5,
    ),
    # the uneven workload: the 2,000-protein database, variable Met oxidation
    # and S/T/Y phosphorylation with at most 3 sites (modification vectors
    # summing below 4), 50,000 spectra at 1+..4+, half of them from modified
    # peptides; per-spectrum modified-candidate counts are heavy-tailed
    "mods-2k-50k": dict(
        proteins=2000, spectra=50000, seed=5, missed=1, p_rule=1, min_len=7, max_len=30,
        charges=(1, 2, 3, 4), charge_p=(0.1, 0.4, 0.35, 0.15), static={}, double3=True,
        topn=5, var={"M": 15.9949, "S": 79.9663, "T": 79.9663, "Y": 79.9663}, max_ma=4,
    ),
}
