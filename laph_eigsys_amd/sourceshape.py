"""Source shape of a stored eigensystem: the host program of sourceshape.cpp
with sourceshape_funcs.cpp:31-136 and the writers of read_write.cpp:206-235.

    python -m laph_eigsys_amd.sourceshape CONFIG TIMESLICE NEV [--params parameters.txt]

Reads <in_path>/eigenvectors.%04d.%03d (read_evectors_bin_ts, with its
trace guard), computes on the device psi(r) = sum_x |V_{x+r} V_x^H|_F for every
r in {0, 2, 4, ..}^3 below L, averages psi over vectors r of equal length
(avg_radii) and writes

    sourceshape_nev<NEV>.<config>.<timeslice>            "#r<TAB>psi(r)" + "r psi" lines
    sourceshape_nev<NEV>.%04d.%03d.bin                   (r, psi) double pairs

into the working directory, as the reference does.
"""
import argparse
import math
import struct
import sys

from . import lapheig as lp
from .ev_ts import IO


def results(L, psi):
    """[((r1, r2, r3), |r|, psi)] in the order of the reference's loops."""
    half = (L + 1) // 2
    out = []
    i = 0
    for r1 in range(0, L, 2):
        for r2 in range(0, L, 2):
            for r3 in range(0, L, 2):
                out.append(((r1, r2, r3), math.sqrt(r1 * r1 + r2 * r2 + r3 * r3), float(psi[i])))
                i += 1
    assert i == half ** 3
    return out


def avg_radii(res):
    """sourceshape_funcs.cpp:102-136: psi averaged over radii equal within
    10e-12, in order of first appearance -> [(r, psi_avg)]."""
    acc = []
    for _, rad, val in res:
        hit = False
        for a in acc:
            if abs(rad - a[1]) < 10e-12:
                a[0] += 1
                a[2] += val
                hit = True
        if not hit:
            acc.append([1, rad, val])
    return [(a[1], a[2] / a[0]) for a in acc]


def write_sourceshape_ascii(prefix, config, tslice, nev, avg):
    """read_write.cpp:206-222: "<prefix>_nev<nev>.<config>.<tslice>" (no zero
    padding), a header line and "r psi" at 12 significant digits."""
    fn = f"{prefix}_nev{nev}.{config}.{tslice}"
    with open(fn, "w") as f:
        f.write("#r\tpsi(r)\n")
        for r, v in avg:
            f.write(f"{r:.12g} {v:.12g}\n")
    return fn


def write_sourceshape_bin(prefix, config, tslice, nev, avg):
    """read_write.cpp:225-235: "<prefix>_nev%d.%04d.%03d.bin", (r, psi) doubles."""
    fn = f"{prefix}_nev{nev}.{config:04d}.{tslice:03d}.bin"
    with open(fn, "wb") as f:
        for r, v in avg:
            f.write(struct.pack("<dd", r, v))
    return fn


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("config", type=int)
    ap.add_argument("timeslice", type=int)
    ap.add_argument("nev", type=int)
    ap.add_argument("--params", default="parameters.txt")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args(argv)
    pars = IO().set_values(args.params)
    L = pars.get_int("LX")
    V = lp.read_evectors_bin_ts(pars.get_path("in"), "eigenvectors", args.config, args.timeslice,
                                pars.get_int("MAT_ENTRIES"), args.nev)
    avg = avg_radii(results(L, lp.source_shape(L, V, args.device)))
    for r, v in avg:
        print(r, v)
    write_sourceshape_ascii("sourceshape", args.config, args.timeslice, args.nev, avg)
    write_sourceshape_bin("sourceshape", args.config, args.timeslice, args.nev, avg)
    return 0


if __name__ == "__main__":
    sys.exit(main())
