"""Helpers for running the ms10 driver in tests."""
import os
import struct
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MS10 = os.path.join(ROOT, "wloop_amd", "ms10")

WLOOP = """[Wloop]
msl          {msl}
nss          {nss}
mwlt         {mwlt}
mwlr         {mwlr}
alpha_action 1.0 1.0 0.5
proj_iter    4
op_smear     APE
alpha        0.5
"""


def write_input(path, d, name="run", first=1, last=2, step=1, bc=3, lattice=(8, 4, 4, 4),
                nss=(0, 1, 2), mwlt=3, mwlr=2, synthetic=(7, 0.4), precision=0, cnfg_dir=None):
    os.makedirs(os.path.join(d, "log"), exist_ok=True)
    os.makedirs(os.path.join(d, "dat"), exist_ok=True)
    s = f"[Run name]\nname         {name}\n\n[Directories]\nlog_dir      {d}/log\ndat_dir      {d}/dat\n"
    if cnfg_dir:
        s += f"cnfg_dir     {cnfg_dir}\n"
    s += f"\n[Configurations]\nfirst        {first}\nlast         {last}\nstep         {step}\n\n"
    s += f"[Boundary conditions]\ntype         {bc}\n\n"
    s += WLOOP.format(msl=len(nss), nss=" ".join(map(str, nss)), mwlt=mwlt, mwlr=mwlr)
    if lattice:
        s += "\n[Lattice]\nsize {} {} {} {}\n".format(*lattice)
    if synthetic:
        s += f"\n[Synthetic]\nseed {synthetic[0]}\neps  {synthetic[1]}\n"
    s += f"\n[Device]\nprecision {precision}\ndevice 0\n"
    with open(path, "w") as f:
        f.write(s)


def run(inp, cwd, *extra, timeout=300):
    return subprocess.run([MS10, "-i", inp, *extra], cwd=cwd, capture_output=True, text=True,
                          timeout=timeout)


def read_dat(path, wls):
    with open(path, "rb") as f:
        head = struct.unpack("<5i", f.read(20))
        recs = []
        while True:
            b = f.read(4)
            if len(b) < 4:
                break
            nc = struct.unpack("<i", b)[0]
            recs.append((nc, np.frombuffer(f.read(8 * wls), dtype="<f8").copy()))
    return head, recs
