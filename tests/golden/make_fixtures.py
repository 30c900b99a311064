"""Write the golden fixtures of the HD+H J=0 example of the reference.

Reads the reference's own run log and probability table
(EXAMPLES/HD+H_collision/J000-check/sal.000 and S2prod.v00.J000.k00008 under
the reference tree) and stores the numbers they hold as data:
  hd_h_j000_pins.json      scalars printed during set-up
  hd_h_j000_angles.txt     the 140 Gauss-Legendre angles (degrees)
  hd_h_j000_norms.txt      |psi_k|^2 for k = 1..4001 (check)
  hd_h_j000_S2prod.txt     the 500-energy probability table
  hd_h_j000_reac_levels.txt  reactant (0-1) levels: ielec j v E - eref (eV)
  hd_h_j000_prod_levels.txt  product (0-2) levels: j v E - eref (eV)
  hd_h_input.dat           the example's input.dat
Run only where the reference tree exists; the outputs are committed.
"""
import json
import os
import re
import sys

REF = os.environ.get("MW3_REFERENCE", "/root/reference")
D = os.path.join(REF, "EXAMPLES", "HD+H_collision", "J000-check")
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    lines = open(os.path.join(D, "sal.000")).read().splitlines()
    pins = {}

    def grab(pattern, names, conv=float):
        for ln in lines:
            m = re.search(pattern, ln)
            if m:
                for n, g in zip(names, m.groups()):
                    pins[n] = conv(g)
                return
        raise ValueError("pattern not found: " + pattern)

    num = r"([-+0-9.Ee]+)"
    grab(r"Absorption starts at ir1,ir2 =\s+(\d+)\s+(\d+)", ["ir1abs", "ir2abs"], int)
    grab(r"E_kin\^1\(max\)=\s+" + num + r"\s+E_kin\^2\(max\)=\s+" + num, ["ekin1_max_eV", "ekin2_max_eV"])
    grab(r"E_kin\^1\(min\)=\s+" + num + r"\s+E_kin\^2\(min\)=\s+" + num, ["ekin1_min_eV", "ekin2_min_eV"])
    grab(r"r20 =\s+" + num + r"\s+k=\s+" + num + r"\s+alpha=\s+" + num, ["r20", "xk0", "alpha0"])
    grab(r"norm of incident gaussian in R2=\s+" + num, ["norm_gauss_R2"])
    grab(r"total flux evaluated at index ir4flux=\s+(\d+)", ["ir4flux"], int)
    grab(r"Norm ini. wvp in energy=\s+" + num, ["norm_energy"])
    grab(r"Emin=\s+" + num + r"\s+Emax=\s+" + num, ["emin_eV", "emax_eV"])
    grab(r"Norm of initial wavepacket=\s+" + num, ["norm_initial"])
    grab(r"Mean energy \(eV\) and norm of initial wvp=\s+" + num, ["emean_eV"])
    # product-coordinate sampling range printed by changeini (iprod = 2)
    grab(r"rpbmin,rpbmax=\s+" + num + r"\s+" + num, ["rpbmin", "rpbmax"])
    json.dump(pins, open(os.path.join(HERE, "hd_h_j000_pins.json"), "w"), indent=1, sort_keys=True)

    # angles: the block printed after "Angular grid to plot pes and wvp"
    i0 = next(i for i, l in enumerate(lines) if "Angular grid to plot pes and wvp" in l)
    ang = []
    for l in lines[i0 + 1:]:
        t = l.split()
        if len(t) == 2 and t[0].isdigit():
            ang.append((int(t[0]), t[1]))
        elif ang and "norealproc" in l:
            break
    with open(os.path.join(HERE, "hd_h_j000_angles.txt"), "w") as f:
        for i, a in ang:
            f.write("%d %s\n" % (i, a))

    # norm series
    i0 = next(i for i, l in enumerate(lines) if re.match(r"\s+k\s+norm\s*$", l))
    norms = []
    for l in lines[i0 + 1:]:
        t = l.split()
        if len(t) == 2 and t[0].isdigit():
            norms.append((int(t[0]), t[1]))
    with open(os.path.join(HERE, "hd_h_j000_norms.txt"), "w") as f:
        for k, v in norms:
            f.write("%d %s\n" % (k, v))

    # the example's input file (the namelists of the run the log belongs to)
    src_in = os.path.join(REF, "EXAMPLES", "HD+H_collision", "input.dat")
    with open(src_in) as f, open(os.path.join(HERE, "hd_h_input.dat"), "w") as g:
        g.write(f.read())

    # diatomic levels, as radial_functions01_read and product_radialf_read
    # print them after subtracting the reference energy
    def levels(header, ncols, name):
        i0 = next(i for i, l in enumerate(lines) if re.match(header, l))
        rows = []
        for l in lines[i0 + 2:]:
            t = l.split()
            if not t and not rows:
                continue
            if len(t) == ncols and all(re.match(r"^-?\d+$", x) for x in t[:-1]):
                rows.append(t)
            else:
                break
        with open(os.path.join(HERE, name), "w") as f:
            for t in rows:
                f.write(" ".join(t) + "\n")
        return len(rows)

    nr = levels(r"\s*ie\s+j\s+v\s+E\(eV\)", 4, "hd_h_j000_reac_levels.txt")
    npr = levels(r"\s*j\s+v\s+E\(eV\)", 3, "hd_h_j000_prod_levels.txt")
    print("levels:", nr, npr)

    src = os.path.join(D, "S2prod.v00.J000.k00008")
    with open(src) as f, open(os.path.join(HERE, "hd_h_j000_S2prod.txt"), "w") as g:
        g.write(f.read())
    print("pins:", len(pins), "angles:", len(ang), "norms:", len(norms))


if __name__ == "__main__":
    sys.exit(main())
