"""Writers for the reference's data_in case folders (two-dim and one-dim).

A case folder holds config.txt ("index value" lines) and RHO/U/V/P.txt: one
text line per grid line in y (bottom first), n_x numbers per line.  The
reader (src/file_io/io_control.c:137-142,194-200,257) accepts only the
characters 0-9 . - e E, so numbers are written with 17 significant digits and
never with a '+' in the exponent.

    python -m hydrogrp.datain four_quadrant 1024 1024 <dir>
"""
import os
import sys

import numpy as np


def fmt(x):
    """Round-trippable repr of a float without '+' (strtod reads it back exactly)."""
    s = "%.17g" % float(x)
    return s.replace("e+", "e")


def write_table(path, field):
    """field[j, i] (x index j, y index i) -> n_y lines of n_x numbers."""
    field = np.asarray(field, dtype=np.float64)
    with open(path, "w") as f:
        for i in range(field.shape[1]):
            f.write("\t".join(fmt(v) for v in field[:, i]) + "\n")


def write_case_2d(folder, init, config_lines, time_plot=None):
    os.makedirs(folder, exist_ok=True)
    with open(os.path.join(folder, "config.txt"), "w") as f:
        for idx, val in config_lines:
            f.write(f"{idx}\t{fmt(val)}\n")
    for name in ("RHO", "U", "V", "P"):
        write_table(os.path.join(folder, name + ".txt"), init[name.lower()])
    if time_plot is not None:
        with open(os.path.join(folder, "time_plot.txt"), "w") as f:
            for t in time_plot:
                f.write(fmt(t) + "\n")


def four_quadrant(m, n):
    """2D Riemann problem of data_in/two-dim/RP2D_Positive/Config3 on an m x n grid."""
    out = {k: np.empty((m, n)) for k in ("rho", "u", "v", "p")}
    west = (np.arange(m) + 0.5) / m < 0.5
    south = (np.arange(n) + 0.5) / n < 0.5
    states = {(True, True): (0.138, 1.206, 1.206, 0.029), (True, False): (0.5323, 1.206, 0.0, 0.3),
              (False, True): (0.5323, 0.0, 1.206, 0.3), (False, False): (1.5, 0.0, 0.0, 1.5)}
    for (w, s), st in states.items():
        mask = np.outer(west == w, south == s)
        for k, val in zip(("rho", "u", "v", "p"), st):
            out[k][mask] = val
    return out


def main(argv):
    if len(argv) != 5 or argv[1] != "four_quadrant":
        print(__doc__)
        return 2
    m, n, folder = int(argv[2]), int(argv[3]), argv[4]
    cfg = [(1, 0.3), (4, 1e-9), (6, 1.4), (7, 0.45), (10, 1.0 / m), (11, 1.0 / n), (17, -4), (18, -4), (33, 1)]
    write_case_2d(folder, four_quadrant(m, n), cfg)
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
