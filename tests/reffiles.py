"""Writers of the reference's input files for tests (io.cpp:112-259 formats)."""
import os

import numpy as np


def write_reference_files(tmp, sysd):
    """Writes bonds/angles/vdw/init in the reference formats (io.cpp:112-259)."""
    n = sysd.n
    with open(os.path.join(tmp, "bonds.dat"), "w") as f:
        f.write(f"nbnd {n - 1}\n")
        for i in range(n - 1):
            f.write(f"{i + 1} {i + 2} {sysd.bond_r0[i]:.6f}\n")
    with open(os.path.join(tmp, "angles.dat"), "w") as f:
        f.write(f"nang {n - 2}\n")
        for i in range(n - 2):
            f.write(f"{i + 1} {i + 2} {i + 3} 0.0\n")
    nat = {(int(a), int(b)): float(r) for a, b, r in zip(sysd.nat_i, sysd.nat_j, sysd.nat_r0)}
    x = sysd.xyz.astype(np.float64)
    lines = []
    for i in range(n):
        for j in range(i + 3, n):
            r = nat.get((i, j), max(8.5, float(np.linalg.norm(x[i] - x[j]))))
            lines.append(f"{i + 1} {j + 1} {r:.6f} 1 1\n")
    with open(os.path.join(tmp, "vdw.dat"), "w") as f:
        f.write(f"ncon_att {len(nat)} ncon_rep {len(lines) - len(nat)}\n")
        f.writelines(lines)
    with open(os.path.join(tmp, "init.dat"), "w") as f:
        f.write(f"nbead {n}\n")
        for i in range(n):
            f.write(f"{i + 1} {x[i, 0]:.6f} {x[i, 1]:.6f} {x[i, 2]:.6f}\n")


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOP_DIR = os.path.join(ROOT, "sop_amd")


def build_shim_driver(out_dir):
    """Compiles tests/helpers/refshim_driver.c against libsop_refshim.so."""
    import subprocess
    exe = os.path.join(str(out_dir), "refshim_driver")
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-o", exe,
                    os.path.join(ROOT, "tests", "helpers", "refshim_driver.c"), "-L" + SOP_DIR,
                    "-lsop_refshim", "-lsop", "-Wl,-rpath," + SOP_DIR, "-lm"], check=True)
    return exe


def write_command_file(path, lines):
    """A command file in the reference's grammar (io.cpp:9-110)."""
    with open(path, "w") as f:
        for ln in lines:
            f.write(ln.rstrip(";") + ";\n")
