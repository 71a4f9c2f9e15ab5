"""The reading rules of the reference's post-processing, restated.

post/fileTreat.py (kreeyto/MULTIC-BUBBLE-CUDA) is the consumer the driver's
output must satisfy. This module restates its rules (it is not imported):
  - getMacrSteps (fileTreat.py:23-34): files matching "*phi*.bin" in the
    folder; the step is the text after the last "phi", minus ".bin", as int.
  - getSimInfo (fileTreat.py:37-124): the first line of "*info*.txt" that
    contains a key substring; the value is the text after the last space.
  - readFileMacr3D (fileTreat.py:127-149): raw values of the info file's
    Precision ('f' for float, 'd' for double), reshaped to (NZ_TOTAL, NY, NX)
    in C order, then axes 0 and 2 swapped, giving [x, y, z].
"""
import glob
import os

import numpy as np

KEYS = [("ID", "Simulation ID", str), ("Prc", "Precision", str), ("NX", "NX", int), ("NY", "NY", int),
        ("NZ", "NZ:", int), ("NZ_TOTAL", "NZ_TOTAL", int), ("Tau", "Tau", float), ("Umax", "Umax", float),
        ("Nsteps", "Nsteps", int), ("FX", "FX", float), ("FY", "FY", float), ("FZ", "FZ", float)]


def macr_steps(path, name="phi"):
    files = sorted(glob.glob(path + "*" + name + "*.bin"))
    return sorted({int(f.split(name)[-1][:-4]) for f in files})


def sim_info(path):
    fname = glob.glob(path + "*info*.txt")[0]
    with open(fname) as fh:
        lines = [l.strip() for l in fh.readlines()]
    info = {}
    for key, sub, conv in KEYS:
        hits = [conv(t.split(" ")[-1]) for t in lines if sub in t]
        if hits:
            info[key] = hits[0]
    return info


def read_macr(fname, path):
    info = sim_info(path)
    prc = {"double": "d", "float": "f"}[info["Prc"]]
    vec = np.fromfile(fname, prc)
    return np.swapaxes(vec.reshape((info["NZ_TOTAL"], info["NY"], info["NX"]), order="C"), 0, 2)


def macr_file(path, sim_id, name, step):
    return os.path.join(path, f"{sim_id}_{name}{step:06d}.bin")
