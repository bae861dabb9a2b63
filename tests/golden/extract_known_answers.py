"""Writes known_answers.json from the reference's saved outputs.

Reads (as text) psi3/output.c1_fzc and psi4/hof/output.{c1_nofzc,cs_fzc}
under the reference tree and records the printed energies, orbital
energies, the CCSD iteration trace, the largest amplitudes and the (T)
triplet counts.  Run once in the build container:
    python tests/golden/extract_known_answers.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def floats(s):
    return [float(x) for x in re.findall(r"-?\d+\.\d+(?:[eE][-+]?\d+)?", s)]


def psi3(path):
    txt = open(path).read().splitlines()
    out = {"source": "psi3/output.c1_fzc"}
    for i, l in enumerate(txt):
        if "Nuclear Repulsion Energy    =" in l:
            out["e_nuc"] = floats(l)[-1]
        elif "Frozen Core Energy          =" in l:
            out["e_fzc"] = floats(l)[-1]
        elif "Total SCF Energy (chkpt)    =" in l:
            out["e_scf"] = floats(l)[-1]
        elif "lowest  eigenvalue of the overlap" in l:
            out["s_min_eig"] = float(l.split()[-1])
        elif "highest eigenvalue of the overlap" in l:
            out["s_max_eig"] = float(l.split()[-1])
        elif "CCSD Energy    =" in l:
            out["e_ccsd_total"] = floats(l)[-1]
        elif "CCSD(T) Energy =" in l:
            out["e_ccsd_t_total"] = floats(l)[-1]
        elif l.strip().startswith("Doubly occupied orbitals"):
            vals = []
            j = i + 1
            while "Unoccupied" not in txt[j]:
                vals += [float(x) for x in re.findall(r"-?\d+\.\d{6}", txt[j])]
                j += 1
            occ = vals
            vals = []
            j += 1
            while txt[j].strip() and "*" not in txt[j]:
                vals += [float(x) for x in re.findall(r"-?\d+\.\d{6}", txt[j])]
                j += 1
            out["eps_occ"] = sorted(occ)
            out["eps_vir"] = sorted(vals)
        elif "Iter   Correlation Energy" in l:
            tr = []
            j = i + 2
            while txt[j].strip() and not txt[j].strip().startswith("Computation"):
                f = txt[j].split()
                if len(f) >= 2 and f[0].isdigit():
                    tr.append(float(f[1]))
                j += 1
            out["ccsd_trace"] = tr
        elif "Largest T1 Amplitudes" in l:
            out["t1_largest"] = [abs(float(txt[i + 1 + k].split()[-1])) for k in range(20)]
        elif "Largest T2 Amplitudes" in l:
            out["t2_largest"] = [abs(float(txt[i + 1 + k].split()[-1])) for k in range(20)]
    return out


def psi4(path, name):
    txt = open(path).read().splitlines()
    out = {"source": name}
    for l in txt:
        if "Nuclear Repulsion Energy =" in l:
            out["e_nuc"] = floats(l)[-1]
        elif "@RHF Final Energy:" in l:
            out["e_scf"] = floats(l)[-1]
        elif l.strip().startswith("MP2 correlation energy"):
            out["e_mp2"] = floats(l)[-1]
        elif l.strip().startswith("CCSD correlation energy"):
            out["e_ccsd"] = floats(l)[-1]
        elif l.strip().startswith("(T) energy"):
            out["e_t"] = floats(l)[-1]
        elif "Number of ijk index combinations" in l:
            out["n_triplets"] = int(l.split()[-1])
    # the SCF electronic dipole moment (a.u., about the origin): the line
    # after the first "Electronic Dipole Moment: (a.u.)" header
    for k, l in enumerate(txt):
        if "Electronic Dipole Moment: (a.u.)" in l:
            out["scf_electronic_dipole"] = floats(txt[k + 1])[:3]
            break
    return out


def main(ref_root):
    data = {
        "h2o_ccpvdz_fc1": psi3(os.path.join(ref_root, "psi3", "output.c1_fzc")),
        "hof_sto3g_c1_allelec": psi4(os.path.join(ref_root, "psi4", "hof", "output.c1_nofzc"),
                                     "psi4/hof/output.c1_nofzc"),
        "hof_sto3g_fc2": psi4(os.path.join(ref_root, "psi4", "hof", "output.cs_fzc"), "psi4/hof/output.cs_fzc"),
    }
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(data, f, indent=1)
    print(json.dumps({k: sorted(v) for k, v in data.items()}))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
