"""Write the golden fixtures taken from the reference tree (run in a container that has it).

Outputs (committed):
  reference_outputs.npz  -- the reference's shipped results of input_file.txt (96^3, 216 spheres,
                            100 steps): velocity_profile_z.txt (94 x 3 plane means, printed %8.6f)
                            and matlab_postproc/velocity_field_centerline.txt (94 x 94 x-z plane
                            at y = 48, printed %9.7f).  Data only.
  mrt_tables.json        -- the numeric coefficient tables of the reference's moment transform
                            (GPU_program.cl:411-428, rows of M) and of its hard-coded inverse
                            (GPU_program.cl:576-594), parsed from the source text into numbers.
  inputs/*.txt           -- the reference's three input files (run inputs, data).

Usage: python tests/golden/make_reference_fixtures.py [/root/reference]
"""
import json
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def parse_rows(lines, var):
    """Parse 'var[k] = a*n[0] + ...' statements into a 19x19 coefficient table."""
    T = np.zeros((19, 19))
    pat = re.compile(r"^\s*%s\[(\d+)\]\s*=\s*(.*);" % var)
    term = re.compile(r"([+-]?)\s*(?:([0-9.]+(?:E[+-]?\d+)?)f?\*)?%s\[(\d+)\]" % ("n" if var == "mn" else "smn"))
    for ln in lines:
        m = pat.match(ln)
        if not m:
            continue
        k = int(m.group(1))
        rhs = m.group(2)
        if rhs.strip() in ("0.0", "0"):
            continue
        for sgn, coef, q in term.findall(rhs):
            c = float(coef) if coef else 1.0
            T[k, int(q)] += -c if sgn == "-" else c
    return T


def main(ref):
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    prof = np.loadtxt(os.path.join(ref, "velocity_profile_z.txt"))
    with open(os.path.join(ref, "velocity_profile_z.txt")) as f:
        prof_text = f.read()
    cl = np.loadtxt(os.path.join(ref, "matlab_postproc", "velocity_field_centerline.txt"))
    np.savez_compressed(os.path.join(HERE, "reference_outputs.npz"),
                        profile=prof.astype(np.float64), profile_text=np.array(prof_text),
                        centerline_idx=cl[:, :3].astype(np.int16), centerline_u=cl[:, 3:].astype(np.float64))
    with open(os.path.join(ref, "GPU_program.cl")) as f:
        src = f.read().splitlines()
    M = parse_rows(src[409:429], "mn")          # lines 410-429
    Minv = parse_rows(src[575:595], "msmn")     # lines 576-595
    with open(os.path.join(HERE, "mrt_tables.json"), "w") as f:
        json.dump({"source": "GPU_program.cl:411-428 (M), 576-594 (inverse)",
                   "M": M.tolist(), "Minv": Minv.tolist()}, f)
    for name in ("input_file.txt", "input_file_power_law.txt", "input_file_casson.txt"):
        shutil.copy(os.path.join(ref, name), os.path.join(HERE, "inputs", name))
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
