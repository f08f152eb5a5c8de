"""Writes tests/golden/traj_spiral_R4.npz from the reference's data file
data_traj_spiral_R4.mat (a MAT-v5 file read with scipy.io.loadmat, which
executes nothing from the file).  The arrays are stored unchanged:
phs_spha [15000 x 16] and phs_conc [15000 x 4] (rad/m^order, 1 us dwell),
datatime [1 x 5268] (us) and tdwelltraj (us).  Run once in a container that
has the reference checked out:

    python tests/golden/make_traj_fixture.py /root/reference/data_traj_spiral_R4.mat
"""
import os
import sys

import numpy as np
import scipy.io as sio


def main(src):
    d = sio.loadmat(src)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "traj_spiral_R4.npz")
    np.savez_compressed(out, phs_spha=d["phs_spha"].astype(np.float64), phs_conc=d["phs_conc"].astype(np.float64),
                        datatime=d["datatime"].astype(np.float64),
                        tdwelltraj=np.asarray(d["tdwelltraj"], dtype=np.float64))
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data_traj_spiral_R4.mat")
