"""Frame output in the reference's Gnuplot binary matrix format
(shear_sim::write_files / shear_sim::output, sb/main/shear_sim.cc:659-726).

File layout (float32): a header row [l, x_0 .. x_{l-1}], then one row per
grid line [y_j, f_0 .. f_{l-1}], where l = m for cell fields (p, q, s, tau,
chi, tem, dev; cell centres) and l = m+1 for node fields (u, v, X, Y, cu;
the periodic column repeated, X shifted by the domain width).  The values
are the FP64 fields rounded to float32, as the reference's fwrite of a
float buffer does.
"""
import os

import numpy as np

# mode -> (file prefix, plane or derived quantity); shear_sim.cc:659-675
MODES = {0: "u", 1: "v", 2: "p", 3: "q", 4: "s", 5: "tau", 6: "chi", 7: "tem", 8: "dev", 9: "X",
         10: "Y", 11: "proj"}
# fields written by write_files (fflags = 4|128|256|512|1024, shear_sim.cc:660)
DEFAULT_FLAGS = 4 | 128 | 256 | 512 | 1024


def field_values(planes, mode, TZ, cu=None):
    """the (n+2, m) array (rows -1..n) that output(mode) reads"""
    if mode in (0, 1, 2, 3, 4, 5, 6, 9, 10):
        plane = {0: 0, 1: 1, 2: 2, 3: 3, 4: 4, 5: 5, 6: 6, 9: 7, 10: 8}[mode]
        return planes[plane]
    if mode == 7:
        return planes[6] * TZ
    if mode == 8:
        q, s, t = planes[3], planes[4], planes[5]
        return np.sqrt(3 * q * q + s * s + t * t)
    if mode == 11:
        if cu is None:
            raise ValueError("mode 11 needs the cu plane")
        return cu
    raise ValueError("unknown output mode %d" % mode)


def gnuplot_matrix(params, planes, mode, TZ, cu=None):
    """bytes of the file shear_sim::output(prefix, mode, sn) writes"""
    m, n = int(params["m"]), int(params["n"])
    ax, bx, ay, by = (float(params[k]) for k in ("ax", "bx", "ay", "by"))
    dx, dy = (bx - ax) / m, (by - ay) / n
    st = 2 <= mode <= 8
    l = m if st else m + 1
    vals = field_values(planes, mode, TZ, cu)
    rows = n if st else n + 1
    out = np.empty((rows + 1, l + 1), dtype=np.float32)
    out[0, 0] = l
    i = np.arange(l, dtype=np.float64)
    out[0, 1:] = ax + ((i + 0.5) * dx if st else i * dx)
    j = np.arange(rows, dtype=np.float64)
    out[1:, 0] = ay + ((j + 0.5) * dy if st else j * dy)
    out[1:, 1:m + 1] = vals[1:rows + 1, :]
    if not st:
        first = out[1:, 1].astype(np.float64)
        out[1:, m + 1] = (first + bx - ax) if mode == 9 else first
    return out.tobytes()


def write_field(directory, prefix, params, planes, mode, sn, TZ, cu=None):
    path = os.path.join(directory, "%s.%d" % (prefix, sn))
    with open(path, "wb") as f:
        f.write(gnuplot_matrix(params, planes, mode, TZ, cu))
    return path


def write_files(directory, params, planes, k, TZ, flags=DEFAULT_FLAGS, cu=None):
    """shear_sim::write_files(k): every field selected by flags (bit = mode)"""
    return [write_field(directory, MODES[mode], params, planes, mode, k, TZ, cu)
            for mode in range(12) if flags & (1 << mode)]
